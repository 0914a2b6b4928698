// MD step kernels of the EI production loop (src/md.f90:285-682, it > 0,
// method < 3, icoll = 0), one 64-lane workgroup per trajectory.
//
//   k_md_pre : step counters, temperature and energy averages, drift/SCF
//              error exits, charge averaging, leap-frog (md.f90:287-394)
//   (egrad   : k_setup -> k_scc -> k_grad on the trajectories still running)
//   k_md_post: checkqc, heating by impactscale, setetemp, fragment
//              connectivity, fragment masses/temperatures, fragmentation
//              exits (md.f90:396-679)
//
// Floating-point bookkeeping reproduces the Fortran operation order exactly:
// FMA contraction is off from here on, sums that the reference accumulates
// sequentially are accumulated sequentially by lane 0, and the reference's
// default-real (REAL(4)) literals are float constants widened to double.
#pragma once
#include "k_grad.h"

#pragma clang fp contract(off)

namespace qx {

constexpr double kKB = 3.166808578545117e-06;   // xtb_mctc_constants.f90:35
constexpr double kFsToAu = 41.3413733365614;    // xtb_mctc_convert.f90:57

enum MdExit { kRunning = 0, kErrScf = 1, kErrDrift = 2, kNfrag6 = 3, kNfragExit = 4,
              kFconst = 5, kMorestep = 6, kMaxStep = 7, kHeatFail = 8 };

struct MdTraj {
    int nstep, nfrag, fconst, morestep, kdump, nadd, done, mdok, fragstate, exit_reason;
    int isec;  // fragmentation generation (main.F90:2228): 1 for the first MD of a trajectory
    int max_steps;  // nmax of this generation (main.F90:2296-2320)
    double ekin, epot, tav, epav, ekav, edum, eerror, ttime, dtime, atlast, ekinstart, etemp,
        eimp, fadd, tadd;
    double fragm[10], fragT[10];
    // CID mode (md.f90 with method 3 after a collision, icoll >= 1): cid
    // selects it; count_average is the reference's SAVE flag, handed in and
    // back; check_fragmented, cnt, save_natf the averaged-structure count;
    // add_steps the extension after a new fragment count (cid.f90:232-234)
    int cid, check_fragmented, count_average, cnt, add_steps;
    int save_natf[10];
    // the ion's centre-of-mass speed (m/s, handed in by cid() and back),
    // the temperature without the flight, the ion's mass and last centre
    double new_velo, new_temp, summass, old_cm[3];
};

struct MdParams {
    double tstep, etempin, ax, ieetemp;
    int nfragexit, max_steps;
};

struct MdPtrs {
    MdTraj* traj;
    double* velo;    // 3 per atom
    const double* velof;
    const double* mass;
    const int* imass;
    const int* num;  // atomic numbers
    int* list;       // fragment labels per atom
    double* avchrg;  // running charge sums (md.f90:355-361)
    double* avxyz;   // running coordinate sums, 3 per atom (md.f90:357,363)
    int* active;     // per trajectory: 1 while running
    double* avxyz2;  // CID mode: coordinate sums since the count started, 3 per atom
    double* store;   // CID mode: the averaged structure (store_avxyz), 3 per atom
};

constexpr double kMsToAu = 1.0 / 2.18769126364e+06;  // m/s in au (xtb_mctc_convert.f90:63)
constexpr int kCntSteps = 50;                        // md.f90 cnt_steps

// cid.f90:232-234
QX_DEV int cid_add_steps(int nuc)
{
    int a = 0;
    if (nuc > 10) a = (nuc / 10) * 500;
    if (nuc >= 40) a = (nuc / 10) * 1000;
    return a;
}

// utility.f90:541-562 (sequential sums) of nat atoms
QX_DEV void md_center_of_mass(int nat, const double* mass, const double* x, double cm[3])
{
    double tm = 0.0;
    cm[0] = cm[1] = cm[2] = 0.0;
    for (int i = 0; i < nat; ++i) {
        cm[0] = cm[0] + mass[i] * x[3 * i];
        cm[1] = cm[1] + mass[i] * x[3 * i + 1];
        cm[2] = cm[2] + mass[i] * x[3 * i + 2];
        tm = tm + mass[i];
    }
    cm[0] = cm[0] / tm;
    cm[1] = cm[1] / tm;
    cm[2] = cm[2] / tm;
}

// utility.f90:69-86
QX_DEV double setetemp(int nfrag, double eimp, double ax, double ieetemp)
{
    double etemp = (double)5000.0f + (double)20000.0f * ax;
    if (eimp > 0 && nfrag <= 1) {
        const double tmp = eimp > 0.0 ? eimp : 0.0;
        etemp = etemp + tmp * ieetemp;
    }
    return etemp;
}

// iniqm.f90:684-732 (gnorm counts grad(2,i) twice and never z)
QX_DEV int checkqc(int nat, double& e, const double* g, const double* q, int mchrg)
{
    if (fabs(e) < 1.0e-8) return 0;
    double gn = 0.0;
    for (int i = 0; i < nat; ++i)
        gn = gn + g[3 * i] * g[3 * i] + g[3 * i + 1] * g[3 * i + 1] + g[3 * i + 1] * g[3 * i + 1];
    gn = sqrt(gn);
    if (gn < 1.0e-8 || gn > (double)20.0f) {
        e = 0.0;
        return 0;
    }
    if (mchrg > 0) {
        double mx = -1.0e300;
        for (int i = 0; i < nat; ++i) mx = fmax(mx, q[i]);
        if (fabs(mx) < 1.0e-5) {
            e = 0.0;
            return 0;
        }
    }
    return 1;
}

// sequential sum of per-atom kinetic terms m (vx^2 + vy^2 + vz^2), ekinet order
QX_DEV double ekin_terms_sum(const double* t, int nat)
{
    double e = 0.0;
    for (int i = 0; i < nat; ++i) e = e + t[i];
    return e;
}


// leap-frog for one system by a 64-lane block; returns the kinetic energy of
// the averaged velocities, accumulated atom-major as md.f90:760-768
QX_DEV double dev_leapfrog(int nat, const double* mass, double dt, const double* g, double* x,
                           double* v, double* term)
{
    const int lane = threadIdx.x;
    for (int i = lane; i < nat; i += 64) {
        const double m = mass[i];
        for (int j = 0; j < 3; ++j) {
            const double vold = v[3 * i + j];
            v[3 * i + j] = v[3 * i + j] - (dt * g[3 * i + j] / m);
            const double vavg = 0.5 * (vold + v[3 * i + j]);
            x[3 * i + j] = x[3 * i + j] + (dt * v[3 * i + j]);
            term[3 * i + j] = 0.5 * (m * vavg * vavg);
        }
    }
    __syncthreads();
    __shared__ double s_ke;
    if (lane == 0) {
        double ke = 0.0;
        for (int k = 0; k < 3 * nat; ++k) ke = ke + term[k];
        s_ke = ke;
    }
    __syncthreads();
    return s_ke;
}

// impactscale (impact.f90:12-55); returns the iteration count k (k >= 20000
// is the reference's stop).  The reference raises scal by 0.0002 (a REAL(4)
// literal, a multiple of 2^-36 as a double) per iteration from 0; every
// partial sum below 4 is then exact, so iteration k uses scal = (k - 1) c
// exactly and the iterations are independent.  Each thread evaluates one k
// (the kinetic energy summed in atom order as ekinet does), a pass covers
// blockDim.x consecutive k, and the first k that ends the reference's loop
// is the answer.  Called by all threads of the block (block size a multiple
// of 64, at most QX_BLOCK); atoms: LDS for 5 nat doubles.
QX_DEV int dev_impactscale(int nat, double* velo, const double* mass, const double* velof,
                           double eimp, double ff, double e0, double* atoms)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    __shared__ int s_wmin[QX_BLOCK / 64];
    for (int i = tid; i < nat; i += nt) {
        atoms[5 * i] = velo[3 * i];
        atoms[5 * i + 1] = velo[3 * i + 1];
        atoms[5 * i + 2] = velo[3 * i + 2];
        atoms[5 * i + 3] = mass[i];
        atoms[5 * i + 4] = velof[i];
    }
    __syncthreads();
    const double esoll = eimp * ff + e0;
    const double c = (double)0.0002f;
    constexpr int kNone = 1 << 30;
    int kstop = kNone;
    for (int base = 1; base <= 20000; base += nt) {
        const int k = base + tid;
        int mine = kNone;
        if (k <= 20000) {
            const double scal = (double)(k - 1) * c;
            double e = 0.0;
            for (int i = 0; i < nat; ++i) {
                const double* a = atoms + 5 * i;
                const double f = 1.0 + a[4] * scal;
                const double v0 = a[0] * f, v1 = a[1] * f, v2 = a[2] * f;
                e = e + a[3] * (v0 * v0 + v1 * v1 + v2 * v2);
            }
            e = e * 0.5;
            if (!(esoll - e > (double)0.001f && k < 20000)) mine = k;
        }
        for (int o = 32; o > 0; o >>= 1) {
            const int other = __shfl_xor(mine, o, 64);
            mine = other < mine ? other : mine;
        }
        if ((tid & 63) == 0) s_wmin[tid >> 6] = mine;
        __syncthreads();
        for (int w = 0; w < nt / 64; ++w) kstop = s_wmin[w] < kstop ? s_wmin[w] : kstop;
        __syncthreads();
        if (kstop != kNone) break;
    }
    // the loop leaves scal one increment past the last iteration: k c
    const double scal = (double)kstop * c;
    for (int i = tid; i < nat; i += nt) {
        double* v = velo + 3 * i;
        const double f = 1.0 + atoms[5 * i + 4] * scal;
        v[0] = atoms[5 * i] * f;
        v[1] = atoms[5 * i + 1] * f;
        v[2] = atoms[5 * i + 2] * f;
    }
    __syncthreads();
    return kstop;
}

// fragment_structure(rcut = 3, at1 = 1, at2 = 0) + fragmass masses by a
// 64-lane block (nat <= 64).  Labels are the connected components numbered by
// their lowest atom index, which is what the reference's flood fill yields.
// Returns nfrag; fragm[0..9] receives the masses (amu) of the non-empty labels.
QX_DEV int dev_fragments(int nat, const int* num, const double* x, const double* mass, int* list,
                         double* fragm, unsigned long long* adj)
{
    const int lane = threadIdx.x;
    for (int i = lane; i < nat; i += 64) {
        unsigned long long row = 0ull;
        const double ri = g_rad[num[i] - 1];
        for (int j = 0; j < nat; ++j) {
            if (j == i) continue;
            const int a = i < j ? i : j, b = i < j ? j : i;
            const double dx = x[3 * a] - x[3 * b], dy = x[3 * a + 1] - x[3 * b + 1],
                         dz = x[3 * a + 2] - x[3 * b + 2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            const double rcov = 3.0 * 0.5 * (ri + g_rad[num[j] - 1]);
            if (r < rcov) row |= (1ull << j);
        }
        adj[i] = row;
    }
    __syncthreads();
    __shared__ int s_nfrag;
    if (lane == 0) {
        unsigned long long unassigned = (nat == 64) ? ~0ull : ((1ull << nat) - 1ull);
        int label = 0;
        while (unassigned) {
            ++label;
            unsigned long long comp = unassigned & (~unassigned + 1ull);
            for (;;) {
                unsigned long long grow = comp, c = comp;
                while (c) {
                    const int i = __ffsll((long long)c) - 1;
                    c &= c - 1ull;
                    grow |= adj[i];
                }
                grow &= unassigned;
                if (grow == comp) break;
                comp = grow;
            }
            unsigned long long c = comp;
            while (c) {
                const int i = __ffsll((long long)c) - 1;
                c &= c - 1ull;
                list[i] = label;
            }
            unassigned &= ~comp;
        }
        double fm[10];
        for (int f = 0; f < 10; ++f) fm[f] = 0.0;
        for (int i = 0; i < nat; ++i) {
            const int l = list[i];
            if (l >= 1 && l <= 10) fm[l - 1] = fm[l - 1] + mass[i];
        }
        const double autoamu = (1.0 / 1.660539040e-27) * 9.10938356e-31;
        int nf = 0;
        for (int f = 0; f < 10; ++f) fragm[f] = 0.0;
        for (int f = 0; f < 10; ++f)
            if (fm[f] > 0) {
                fragm[nf] = fm[f] * autoamu;
                ++nf;
            }
        s_nfrag = nf;
    }
    __syncthreads();
    return s_nfrag;
}

__global__ __launch_bounds__(64) void k_md_init(BatchPtrs B, MdPtrs M, MdParams p)
{
    const int sys = blockIdx.x;
    if (!M.active[sys]) return;
    const SysInfo si = B.info[sys];
    const int lane = threadIdx.x, nat = si.nat, off = si.atom_off;
    __shared__ double term[QX_MAX_NAT];
    for (int i = lane; i < nat; i += 64) {
        const double* v = M.velo + 3 * (size_t)(off + i);
        term[i] = M.mass[off + i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    __syncthreads();
    if (lane != 0) return;
    MdTraj& t = M.traj[sys];
    const double ek = ekin_terms_sum(term, nat) * 0.5;  // mdinit.f90:57-76
    t.ekin = ek;
    t.ekinstart = ek;
    t.nadd = (int)((t.tadd + p.tstep) / p.tstep - 1);  // md.f90:265
    t.fadd = p.tstep / (t.tadd + p.tstep);
    t.nfrag = 1;
    t.kdump = 50;
    double e = B.energy[sys];
    checkqc(nat, e, B.grad + 3 * (size_t)off, B.qat + off, si.charge);
    t.epot = e;
    if (e == 0.0) {
        t.done = 1;
        t.exit_reason = kErrScf;
        M.active[sys] = 0;
    }
}

// CID mode start (md.f90:247-255) of the trajectories with t.cid set before
// their first step: the ion's mass and centre, and Ekin without the kinetic
// energy of the flight at the speed new_velo handed over by cid()
__global__ __launch_bounds__(64) void k_md_cid_start(BatchPtrs B, MdPtrs M)
{
    const int sys = blockIdx.x;
    if (threadIdx.x != 0 || !M.active[sys]) return;
    MdTraj& t = M.traj[sys];
    if (!t.cid) return;
    const SysInfo si = B.info[sys];
    const int nat = si.nat, off = si.atom_off;
    md_center_of_mass(nat, M.mass + off, B.xyz + 3 * (size_t)off, t.old_cm);
    double summass = 0.0;
    for (int i = 0; i < nat; ++i) summass = summass + M.mass[off + i];
    t.summass = summass;
    const double vv = t.new_velo * kMsToAu;
    const double e_kin_diff = t.ekin - 0.5 * summass * (vv * vv);
    t.new_temp = (2.0 * e_kin_diff) / (3.0 * kKB * nat);
    t.ekin = e_kin_diff;
    t.add_steps = cid_add_steps(nat);
    t.check_fragmented = 1;
    t.cnt = 0;
}

// Re-layout of a running ensemble: system s of the new batch (Bn, Mn) takes
// the whole state of system src[s] of the old one (Bo, Mo), the same
// molecule at another place: coordinates, gradient, charges, velocities,
// heating factors, masses, element data, fragment labels, running averages,
// the MD record, the electronic temperature, and the SCC's warm start
// (moments of the last five steps, orthogonal basis U, warm flag; the
// Broyden history when keep_hist).  src[s] < 0: a new system, left alone.
__global__ __launch_bounds__(64) void k_md_repack(BatchPtrs Bo, MdPtrs Mo, BatchPtrs Bn, MdPtrs Mn, const int* src,
                                                  int keep_hist, int mem)
{
    const int s = blockIdx.x, o = src[s];
    if (o < 0) return;
    const SysInfo sn = Bn.info[s], so = Bo.info[o];
    const int lane = threadIdx.x, nat = sn.nat;
    const size_t an = sn.atom_off, ao = so.atom_off;
    for (int i = lane; i < nat; i += 64) {
        Bn.qat[an + i] = Bo.qat[ao + i];
        ((double*)Mn.velof)[an + i] = Mo.velof[ao + i];
        ((double*)Mn.mass)[an + i] = Mo.mass[ao + i];
        ((int*)Mn.imass)[an + i] = Mo.imass[ao + i];
        ((int*)Mn.num)[an + i] = Mo.num[ao + i];
        Mn.list[an + i] = Mo.list[ao + i];
        Mn.avchrg[an + i] = Mo.avchrg[ao + i];
    }
    for (int i = lane; i < 3 * nat; i += 64) {
        Bn.xyz[3 * an + i] = Bo.xyz[3 * ao + i];
        Bn.grad[3 * an + i] = Bo.grad[3 * ao + i];
        Mn.velo[3 * an + i] = Mo.velo[3 * ao + i];
        Mn.avxyz[3 * an + i] = Mo.avxyz[3 * ao + i];
        Mn.avxyz2[3 * an + i] = Mo.avxyz2[3 * ao + i];
        Mn.store[3 * an + i] = Mo.store[3 * ao + i];
    }
    for (int i = lane; i < sn.nx; i += 64) {
        Bn.x[sn.x_off + i] = Bo.x[so.x_off + i];
        Bn.xo[sn.x_off + i] = Bo.xo[so.x_off + i];
        Bn.xprev[sn.x_off + i] = Bo.xprev[so.x_off + i];
        Bn.xprev2[sn.x_off + i] = Bo.xprev2[so.x_off + i];
        Bn.xprev3[sn.x_off + i] = Bo.xprev3[so.x_off + i];
        Bn.xprev4[sn.x_off + i] = Bo.xprev4[so.x_off + i];
    }
    const int ld2 = sn.ld * sn.ld;
    for (int i = lane; i < ld2; i += 64) Bn.U[sn.mat_off + i] = Bo.U[so.mat_off + i];
    if (keep_hist) {
        const long long nh = 2LL * mem * sn.nx + 2LL * mem * mem + 2LL * mem + 2LL * sn.nx + 8;
        for (long long i = lane; i < nh; i += 64) Bn.hist[sn.hist_off + i] = Bo.hist[so.hist_off + i];
    }
    if (lane == 0) {
        Mn.traj[s] = Mo.traj[o];
        Bn.warm[s] = Bo.warm[o];
        Bn.iters_total[s] = Bo.iters_total[o];
        Bn.energy[s] = Bo.energy[o];
        ((SysInfo*)Bn.info)[s].kt = so.kt;
    }
}

// LDS of the MD bookkeeping of one trajectory
struct MdLds {
    double term[3 * QX_MAX_NAT];
    double atoms[5 * QX_MAX_NAT];
    unsigned long long adj[QX_MAX_NAT];
    int s_exit, s_more, s_cidop, s_cidcnt;
    double s_T;
};

// CID mode: the averaged structure over the cnt_steps steps after a new
// fragment count (md.f90:496-627).  Thread 0 keeps the count and decides; every
// thread then updates its coordinates of the sums (avxyz2) and the average
// (store_avxyz).  Called by all threads of the block after the connectivity
// (list) of this step.
QX_DEV void md_cid_average(const BatchPtrs& B, const MdPtrs& M, int sys, int nfrag, MdLds& Lm)
{
    enum { kZero = 1, kAdd = 2, kDrop = 4, kFinal = 8 };
    const SysInfo si = B.info[sys];
    const int nat = si.nat, off = si.atom_off;
    MdTraj& t = M.traj[sys];
    if (threadIdx.x == 0) {
        int op = 0, cnt_used = 0;
        if (nfrag > t.check_fragmented) {
            t.count_average = 1;
            t.check_fragmented = nfrag;
            t.max_steps = t.nstep + t.add_steps;
        }
        // a count that falls back drops the average
        if (nfrag < t.check_fragmented && t.count_average) {
            t.cnt = 0;
            op |= kZero;
            t.count_average = 0;
            t.check_fragmented = 1;
        }
        if (t.count_average) {
            t.cnt = t.cnt + 1;
            cnt_used = t.cnt;
            op |= kAdd;
            int natf[10];
            for (int f = 0; f < 10; ++f) natf[f] = 0;
            for (int i = 0; i < nat; ++i) {
                const int l = M.list[off + i];
                if (l >= 1 && l <= 10) natf[l - 1] += 1;
            }
            // a fragment whose atom count changed restarts the count
            for (int f = 0; f < nfrag && f < 10; ++f) {
                if (t.cnt == 1) t.save_natf[f] = natf[f];
                if (natf[f] != t.save_natf[f]) {
                    t.cnt = 0;
                    op |= kDrop;
                    break;
                }
            }
            if (t.cnt == kCntSteps) {
                op |= kFinal;
                t.cnt = 0;
                t.count_average = 0;
            }
        }
        Lm.s_cidop = op;
        Lm.s_cidcnt = cnt_used;
    }
    __syncthreads();
    const int op = Lm.s_cidop;
    if (!op) return;
    const double cnt = (double)Lm.s_cidcnt;
    for (int k = threadIdx.x; k < 3 * nat; k += blockDim.x) {
        const size_t i = 3 * (size_t)off + k;
        double a2 = (op & kZero) ? 0.0 : M.avxyz2[i];
        double st = (op & kZero) ? 0.0 : M.store[i];
        if (op & kAdd) {
            a2 = a2 + B.xyz[i];
            st = a2 / cnt;
        }
        if (op & kDrop) {
            a2 = 0.0;
            st = 0.0;
        }
        if (op & kFinal) a2 = 0.0;  // the average stays in store
        M.avxyz2[i] = a2;
        M.store[i] = st;
    }
}

// one MD loop iteration up to the force call (md.f90:285-395): time,
// averages, drift and SCF checks, running sums, leap-frog.  Any block size
// from 64 threads; lanes from 64 on take no atoms.
QX_DEV void md_pre_system(const BatchPtrs& B, const MdPtrs& M, const MdParams& p, int sys, MdLds& Lm)
{
    if (!M.active[sys]) return;
    const SysInfo si = B.info[sys];
    const int lane = threadIdx.x, nat = si.nat, off = si.atom_off;
    MdTraj& t = M.traj[sys];
    double* term = Lm.term;
    int& s_exit = Lm.s_exit;
    double& s_T = Lm.s_T;
    if (lane == 0) {
        int nstep = t.nstep + 1;
        t.nstep = nstep;
        const double T = t.ekin / ((double)(1.5f * (float)nat) * kKB);
        t.tav = t.tav + T;
        t.epav = t.epav + t.epot;
        t.ekav = t.ekav + t.ekin;
        double eav;
        if (nstep > t.nadd) {
            t.edum = t.edum + t.epot + t.ekin;
            eav = t.edum / (double)(float)(nstep - t.nadd);
        } else {
            eav = t.epot + t.ekin;
        }
        t.eerror = eav - t.epot - t.ekin;
        if (t.nfrag == 1) t.fragT[0] = t.tav / nstep;
        // REAL(4) limits: 0.1 Eh, 0.2 Eh with CID (md.f90:324-328)
        const int err1 = (t.epot == 0.0), err2 = (fabs(t.eerror) > (t.cid ? (double)0.2f : (double)0.1f));
        s_exit = 0;
        if (err1 || err2) {
            t.mdok = (t.nfrag > 1 && t.nfrag <= 4) || t.isec > 1;
            t.exit_reason = err1 ? kErrScf : kErrDrift;
            t.done = 1;
            s_exit = 1;
        } else if (t.kdump > 49) {
            t.kdump = 0;
            t.atlast = 0.0;
        }
        // aTlast sums the temperature without the flight with CID (md.f90:365-366)
        s_T = t.cid ? t.new_temp : T;
    }
    __syncthreads();
    // the flag is cleared only after every wavefront has passed the entry
    // test above, so that all of them take the same barriers
    if (s_exit) {
        if (lane == 0) M.active[sys] = 0;
        return;
    }
    const bool reset = (t.kdump == 0);
    for (int i = lane; i < nat; i += 64) {
        const double prev = reset ? 0.0 : M.avchrg[off + i];
        M.avchrg[off + i] = prev + B.qat[off + i];
        for (int c = 0; c < 3; ++c) {
            const size_t k = 3 * (size_t)(off + i) + c;
            M.avxyz[k] = (reset ? 0.0 : M.avxyz[k]) + B.xyz[k];
        }
    }
    __syncthreads();
    // leap-frog (md.f90:749-773)
    const double ke = dev_leapfrog(nat, M.mass + off, p.tstep, B.grad + 3 * (size_t)off,
                                   B.xyz + 3 * (size_t)off, M.velo + 3 * (size_t)off, term);
    if (lane == 0) {
        t.atlast = t.atlast + s_T;
        t.ekin = ke;
        t.ttime = t.ttime + p.tstep / kFsToAu;
    }
}

// the rest of the iteration after the force call (md.f90:396-682): checkqc,
// heating, connectivity, fragment temperatures, electronic temperature,
// exits.  Any block size from 64 threads.
QX_DEV void md_post_system(const BatchPtrs& B, const MdPtrs& M, const MdParams& p, int sys, MdLds& Lm)
{
    if (!M.active[sys]) return;
    const SysInfo si = B.info[sys];
    const int lane = threadIdx.x, nat = si.nat, off = si.atom_off;
    MdTraj& t = M.traj[sys];
    unsigned long long* adj = Lm.adj;
    int& s_more = Lm.s_more;
    if (lane == 0) {
        double e = B.energy[sys];
        checkqc(nat, e, B.grad + 3 * (size_t)off, B.qat + off, si.charge);
        t.epot = e;
        t.kdump = t.kdump + 1;
        if (t.nfrag == 1) t.morestep = 0;
        if (t.nfrag > 1 && t.dtime < 1.0e-6) t.dtime = t.ttime / (double)1000.0f;
        // heating and setetemp are EI only (md.f90:438-445)
        s_more = (!t.cid && t.nstep <= t.nadd && t.nfrag == 1);
    }
    __syncthreads();
    // heating: impactscale (impact.f90:12-55) with ff = fadd * nstep
    if (s_more) {
        const int k = dev_impactscale(nat, M.velo + 3 * (size_t)off, M.mass + off, M.velof + off,
                                      t.eimp, t.fadd * t.nstep, t.ekinstart, Lm.atoms);
        if (k >= 20000) {
            if (lane == 0) {
                t.done = 1;
                t.exit_reason = kHeatFail;
                M.active[sys] = 0;
            }
            return;
        }
    }
    // connectivity (fragments.f90:93-182) and fragment masses (fragments.f90:10-84)
    double fragm[10];
    const int nfrag = dev_fragments(nat, M.num + off, B.xyz + 3 * (size_t)off, M.mass + off,
                                    M.list + off, fragm, adj);
    // t.cid does not change during a run: the test is uniform over the block
    if (t.cid && nfrag <= 6) md_cid_average(B, M, sys, nfrag, Lm);
    if (lane != 0) return;
    if (!t.cid) {
        // electronic temperature for the next egrad (md.f90:443-445): setetemp
        // runs before fragment_structure, so it sees the previous step's nfrag
        const double dum = t.eimp - t.eimp * (double)(float)t.nstep / (double)(float)t.nadd;
        t.etemp = setetemp(t.nfrag, dum, p.ax, p.ieetemp);
        ((SysInfo*)B.info)[sys].kt = t.etemp * gfn2::kKtoAu;
    }
    for (int f = 0; f < 10; ++f) t.fragm[f] = fragm[f];
    t.nfrag = nfrag;
    // fragment temperatures (md.f90:715-741)
    if (nfrag > 1) {
        double eint[10];
        int cnt[10];
        for (int f = 0; f < 10; ++f) {
            eint[f] = 0.0;
            cnt[f] = 0;
        }
        for (int i = 0; i < nat; ++i) {
            const int j = M.list[off + i];
            if (j < 1 || j > 10) continue;
            const double* v = M.velo + 3 * (size_t)(off + i);
            eint[j - 1] = eint[j - 1] + 0.5 * M.mass[off + i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            cnt[j - 1] = cnt[j - 1] + 1;
        }
        for (int f = 0; f < nfrag && f < 10; ++f)
            t.fragT[f] = eint[f] / ((double)(1.5f * (float)cnt[f]) * kKB);
    }
    // exits (md.f90:458-462, 630-679); later generations exit at two
    // fragments (main.F90:2255)
    const int nfragexit = t.isec > 1 ? 2 : p.nfragexit;
    int reason = kRunning;
    if (nfrag > 6) {
        reason = kNfrag6;
    } else if (t.cid) {
        // the ion's speed from its centre's move over the step, and Ekin
        // without the flight (md.f90:464-487)
        double cm[3];
        md_center_of_mass(nat, M.mass + off, B.xyz + 3 * (size_t)off, cm);
        const double d0 = cm[0] - t.old_cm[0], d1 = cm[1] - t.old_cm[1], d2 = cm[2] - t.old_cm[2];
        const double cm_out = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        t.new_velo = (cm_out / p.tstep) / kMsToAu;
        t.old_cm[0] = cm[0];
        t.old_cm[1] = cm[1];
        t.old_cm[2] = cm[2];
        const double vv = t.new_velo * kMsToAu;
        const double e_kin = 0.5 * t.summass * (vv * vv);
        const double e_kin_diff = t.ekin - e_kin;
        t.new_temp = (2.0 * e_kin_diff) / (3.0 * kKB * nat);
        t.ekin = e_kin_diff;
        // no fragment exits: the run ends at max_steps, which a new fragment
        // count moves to add_steps later (md.f90:504, 672-679)
        if (t.nstep >= t.max_steps) {
            t.fragstate = 1;
            t.mdok = 1;
            reason = kMaxStep;
        }
    } else if (nfrag > nfragexit) {
        t.fragstate = 1;
        t.mdok = 1;
        reason = kNfragExit;
    } else {
        if (nfrag >= 2) t.fconst = t.fconst + 1;
        else t.fconst = 0;
        if (t.fconst > 1000) {
            t.fragstate = 2;
            t.mdok = 1;
            reason = kFconst;
        } else if (nfrag >= nfragexit && (t.morestep = t.morestep + 1) > 250) {
            t.fragstate = 1;
            t.mdok = 1;
            reason = kMorestep;
        } else if (t.nstep >= t.max_steps) {
            t.fragstate = 1;
            t.mdok = 1;
            reason = kMaxStep;
        }
    }
    if (reason != kRunning) {
        t.exit_reason = reason;
        t.done = 1;
        M.active[sys] = 0;
    }
}

__global__ __launch_bounds__(64) void k_md_pre(BatchPtrs B, MdPtrs M, MdParams p)
{
    __shared__ MdLds Lm;
    md_pre_system(B, M, p, blockIdx.x, Lm);
}

// 256 threads: the heating search evaluates 256 increments per pass
__global__ __launch_bounds__(QX_BLOCK) void k_md_post(BatchPtrs B, MdPtrs M, MdParams p)
{
    __shared__ MdLds Lm;
    md_post_system(B, M, p, blockIdx.x, Lm);
}

// Persistent MD: one block of QX_SCC_BLOCK threads per trajectory runs up to
// nsteps whole loop iterations (bookkeeping, setup, SCC, gradient,
// bookkeeping) back to back, so no trajectory waits at a kernel boundary for
// the slowest one.  The phases share one LDS allocation; the dynamic LDS
// holds the setup's Cholesky factor and the SCC's matrices in turn.
union FusedLds {
    SetupLds s;
    SccLds c;
    GradLds g;
    MdLds m;
};

template <bool EIG>
__global__ __launch_bounds__(QX_SCC_BLOCK) void k_md_fused(BatchPtrs B, MdPtrs M, MdParams p, SccOpts opt,
                                                           int nsteps)
{
    static_assert(QX_SCC_BLOCK == QX_BLOCK, "the fused phases assume one block size");
    const int sys = blockIdx.x;
    __shared__ FusedLds L;
    extern __shared__ double dyn[];
#ifdef QX_PHASE_TIMING
    // per-phase cycles of the step (slots 24..28: bookkeeping, setup, SCC,
    // gradient, bookkeeping)
    unsigned long long fz = clock64();
#define QX_FUSED_PH(k)                                                                  \
    do {                                                                                \
        const unsigned long long t_ = clock64();                                        \
        if (threadIdx.x == 0 && sys < kPhaseSys) g_phase[sys * kPhaseSlots + (k)] += t_ - fz; \
        fz = t_;                                                                        \
    } while (0)
#else
#define QX_FUSED_PH(k) \
    do {               \
    } while (0)
#endif
    for (int step = 0; step < nsteps; ++step) {
        if (!M.active[sys]) break;
        md_pre_system(B, M, p, sys, L.m);
        __syncthreads();
        QX_FUSED_PH(24);
        if (!M.active[sys]) break;
        setup_system(B, sys, 1, L.s, dyn);
        __syncthreads();
        QX_FUSED_PH(25);
        scc_system<EIG>(B, opt, sys, L.c, dyn);
        __syncthreads();
        QX_FUSED_PH(26);
        grad_system(B, sys, L.g);
        __syncthreads();
        QX_FUSED_PH(27);
        md_post_system(B, M, p, sys, L.m);
        __syncthreads();
        QX_FUSED_PH(28);
    }
#undef QX_FUSED_PH
}

}  // namespace qx

namespace qx {

// ---- standalone batched drop-ins (one 64-lane block per system) ----
__global__ __launch_bounds__(64) void k_fragments_batch(const int* offsets, const int* num,
                                                        const double* xyz, const double* mass,
                                                        int* list, int* nfrag, double* fragm)
{
    const int sys = blockIdx.x, off = offsets[sys], nat = offsets[sys + 1] - off;
    __shared__ unsigned long long adj[QX_MAX_NAT];
    double fm[10];
    const int nf = dev_fragments(nat, num + off, xyz + 3 * (size_t)off, mass + off, list + off, fm, adj);
    if (threadIdx.x == 0) {
        nfrag[sys] = nf;
        for (int f = 0; f < 10; ++f) fragm[10 * sys + f] = fm[f];
    }
}

__global__ __launch_bounds__(64) void k_leapfrog_batch(const int* offsets, const double* grad,
                                                       const double* mass, double tstep, double* xyz,
                                                       double* velo, double* ekin)
{
    const int sys = blockIdx.x, off = offsets[sys], nat = offsets[sys + 1] - off;
    __shared__ double term[3 * QX_MAX_NAT];
    const double ke = dev_leapfrog(nat, mass + off, tstep, grad + 3 * (size_t)off, xyz + 3 * (size_t)off,
                                   velo + 3 * (size_t)off, term);
    if (threadIdx.x == 0) ekin[sys] = ke;
}

__global__ __launch_bounds__(64) void k_impactscale_batch(const int* offsets, double* velo,
                                                          const double* mass, const double* velof,
                                                          const double* eimp, const double* ff,
                                                          const double* e0, int* iters)
{
    const int sys = blockIdx.x, off = offsets[sys], nat = offsets[sys + 1] - off;
    __shared__ double atoms[5 * QX_MAX_NAT];
    const int k = dev_impactscale(nat, velo + 3 * (size_t)off, mass + off, velof + off, eimp[sys], ff[sys],
                                  e0[sys], atoms);
    if (threadIdx.x == 0) iters[sys] = k;
}

}  // namespace qx
