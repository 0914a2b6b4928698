// Collision MD of QCxMS's CID mode (src/cid.f90:724-1111, one collision of
// cid() after its set-up) for many trajectories, one 64-lane workgroup per
// trajectory in the bookkeeping kernels:
//
//   k_cid_init: the initial checkqc, centre of mass and tracked distance
//               (cid.f90:703, 733-738)
//   k_cid_pre : step counters, average reset, leap-frog of ion and gas
//               (cid.f90:743-770)
//   (egrad    : k_setup -> k_scc -> k_grad of ion + gas, cid.f90:773)
//   k_cid_post: checkqc, the ion's centre-of-mass speed and temperature,
//               connectivity and fragment masses of the ion, the averaged
//               fragment structure, the distance test every dumpdist steps,
//               the collision and fragmentation extensions, the stop test,
//               fragment energies at the end (cid.f90:775-1061)
//
// A system holds the ion's nuc atoms followed by the gas atoms; the last atom
// is the tracked one.  Sums the reference accumulates sequentially are
// accumulated sequentially by lane 0 with FMA contraction off, as in the EI
// MD kernels.
#pragma once
#include "k_md.h"

#pragma clang fp contract(off)

namespace qx {

constexpr double kAutoEvCid = 27.21138505;           // xtb_mctc_convert.f90:27

struct CidTraj {
    // set-up (inputs)
    int nuc, nuc0, ntot, add_steps;
    double tinit, beta, velo_cm, summass;
    // loop state
    int nstep, m, xyzavg_dump, average_dump, distance_dump, step_counter, total_steps,
        check_fragmented, cnt, nfrag, collided, count_average, fragmented, gradfail, done,
        collision_step;
    int save_natf[10];
    double ttime, epot, ekin, old_cm[3], lowest, highest, tav, avgt, atlast, new_velo, ecom,
        velo_diff;
    double fragT[10], eint[10];
};

struct CidParams {
    double tstep;
    int dumpdist, dumpavg, cnt_steps;
};

struct CidPtrs {
    CidTraj* traj;
    double* velo;        // 3 per atom (ion and gas)
    const double* mass;  // per atom
    const int* imass;    // per atom: isotope slot or -1
    const int* num;      // atomic numbers
    int* list;           // fragment labels of the ion's atoms
    double* avxyz;       // coordinate sums since the last reset, 3 per atom
    double* avxyz2;      // sums for the averaged fragment structure
    double* store;       // the averaged fragment structure
    int* active;         // per trajectory: 1 while running
};

// cid.f90:1124-1139: distance of the last atom to cm
QX_DEV double cid_dist(int nuc0, const double* x, const double cm[3])
{
    const double* a = x + 3 * (nuc0 - 1);
    const double d0 = a[0] - cm[0], d1 = a[1] - cm[1], d2 = a[2] - cm[2];
    return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

__global__ __launch_bounds__(64) void k_cid_init(BatchPtrs B, CidPtrs C)
{
    const int sys = blockIdx.x;
    if (threadIdx.x != 0) return;
    const SysInfo si = B.info[sys];
    const int off = si.atom_off;
    CidTraj& t = C.traj[sys];
    double e = B.energy[sys];
    const int ok = B.status[sys] == 0 && checkqc(t.nuc0, e, B.grad + 3 * (size_t)off, B.qat + off, si.charge);
    t.epot = e;
    double cm[3];
    md_center_of_mass(t.nuc, C.mass + off, B.xyz + 3 * (size_t)off, cm);
    for (int c = 0; c < 3; ++c) t.old_cm[c] = cm[c];
    t.lowest = t.highest = cid_dist(t.nuc0, B.xyz + 3 * (size_t)off, cm);
    if (!ok) {
        t.gradfail = 1;
        t.done = 1;
        C.active[sys] = 0;
    }
}

__global__ __launch_bounds__(64) void k_cid_pre(BatchPtrs B, CidPtrs C, CidParams p)
{
    const int sys = blockIdx.x;
    if (!C.active[sys]) return;
    const SysInfo si = B.info[sys];
    const int lane = threadIdx.x, off = si.atom_off;
    CidTraj& t = C.traj[sys];
    __shared__ double term[3 * QX_MAX_NAT];
    __shared__ int s_reset;
    if (lane == 0) {
        t.nstep = t.nstep + 1;
        s_reset = (t.xyzavg_dump == p.dumpavg);
        if (s_reset) t.xyzavg_dump = 0;
        t.ttime = t.ttime + p.tstep * (1.0 / kFsToAu);
        t.distance_dump = t.distance_dump + 1;
        t.xyzavg_dump = t.xyzavg_dump + 1;
        t.average_dump = t.average_dump + 1;
    }
    __syncthreads();
    if (s_reset)
        for (int k = lane; k < 3 * t.nuc; k += 64) C.avxyz[3 * (size_t)off + k] = 0.0;
    const double ke = dev_leapfrog(t.nuc0, C.mass + off, p.tstep, B.grad + 3 * (size_t)off,
                                   B.xyz + 3 * (size_t)off, C.velo + 3 * (size_t)off, term);
    if (lane == 0) t.ekin = ke;
}

__global__ __launch_bounds__(64) void k_cid_post(BatchPtrs B, CidPtrs C, CidParams p)
{
    const int sys = blockIdx.x;
    if (!C.active[sys]) return;
    const SysInfo si = B.info[sys];
    const int lane = threadIdx.x, off = si.atom_off;
    CidTraj& t = C.traj[sys];
    const int nuc = t.nuc;
    const double* x = B.xyz + 3 * (size_t)off;
    const double* mass = C.mass + off;
    double* av = C.avxyz + 3 * (size_t)off;
    double* av2 = C.avxyz2 + 3 * (size_t)off;
    double* st = C.store + 3 * (size_t)off;
    __shared__ unsigned long long adj[QX_MAX_NAT];
    __shared__ int s_stop, s_zero, s_count, s_done;
    __shared__ double s_cnt;
    if (lane == 0) {
        double e = B.energy[sys];
        const int ok = B.status[sys] == 0 &&
                       checkqc(t.nuc0, e, B.grad + 3 * (size_t)off, B.qat + off, si.charge);
        t.epot = e;
        s_stop = !ok;
        if (ok) {
            double cm[3];
            md_center_of_mass(nuc, mass, x, cm);
            const double d0 = cm[0] - t.old_cm[0], d1 = cm[1] - t.old_cm[1], d2 = cm[2] - t.old_cm[2];
            const double cm_out = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            for (int c = 0; c < 3; ++c) t.old_cm[c] = cm[c];
            t.new_velo = (t.nstep != 1) ? (cm_out / p.tstep) / kMsToAu : 0.0;
            // ekinet of the ion (mdinit.f90:57-76)
            const double* v = C.velo + 3 * (size_t)off;
            double ek = 0.0;
            for (int i = 0; i < nuc; ++i)
                ek = ek + mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
            ek = ek * 0.5;
            const double vv = t.new_velo * kMsToAu;
            const double e_velo = 0.5 * t.summass * (vv * vv);
            double new_temp = (2.0 * (ek - e_velo)) / (3.0 * kKB * nuc);
            if (t.nstep == 1) new_temp = t.tinit;
            t.tav = t.tav + new_temp;
            t.m = t.m + 1;
            t.avgt = t.tav / t.m;
            t.ecom = (t.beta * e_velo) * kAutoEvCid;
        }
    }
    __syncthreads();
    if (s_stop) {
        if (lane == 0) {
            t.gradfail = 1;
            t.done = 1;
            C.active[sys] = 0;
        }
        return;
    }
    // connectivity and fragment masses of the ion (cid.f90:839-840)
    double fragm[10];
    const int nfrag = dev_fragments(nuc, C.num + off, x, mass, C.list + off, fragm, adj);
    for (int k = lane; k < 3 * nuc; k += 64) av[k] = av[k] + x[k];
    // the averaged fragment structure (cid.f90:847-958): lane 0 updates the
    // flags and counters, the block the sums
    if (lane == 0) {
        t.nfrag = nfrag;
        s_zero = 0;
        if (nfrag > t.check_fragmented) {
            t.count_average = 1;
            t.check_fragmented = nfrag;
        }
        if (nfrag < t.check_fragmented && t.count_average) {
            t.cnt = 0;
            t.count_average = 0;
            t.check_fragmented = 1;
            s_zero = 1;
        }
        s_count = t.count_average;
        if (s_count) t.cnt = t.cnt + 1;
        s_cnt = (double)t.cnt;
    }
    __syncthreads();
    if (s_zero)
        for (int k = lane; k < 3 * nuc; k += 64) av2[k] = st[k] = 0.0;
    if (s_count)
        for (int k = lane; k < 3 * nuc; k += 64) {
            av2[k] = av2[k] + x[k];
            st[k] = av2[k] / s_cnt;
        }
    __syncthreads();
    if (lane == 0) {
        s_zero = 0;    // a fragment changed its size: the average restarts
        s_done = 0;    // the average of cnt_steps structures is complete
        if (s_count) {
            int natf[10];
            for (int f = 0; f < 10; ++f) natf[f] = 0;
            for (int i = 0; i < nuc; ++i) {
                const int l = C.list[off + i];
                if (l >= 1 && l <= 10) natf[l - 1] += 1;
            }
            for (int f = 0; f < nfrag && f < 10; ++f) {
                if (t.cnt == 1) t.save_natf[f] = natf[f];
                if (natf[f] != t.save_natf[f]) {
                    t.cnt = 0;
                    s_zero = 1;
                    break;
                }
            }
            if (t.cnt == p.cnt_steps) {
                t.cnt = 0;
                t.count_average = 0;
                s_done = 1;
            }
        }
    }
    __syncthreads();
    if (s_zero)
        for (int k = lane; k < 3 * nuc; k += 64) st[k] = av2[k] = 0.0;
    if (s_done)
        for (int k = lane; k < 3 * nuc; k += 64) av2[k] = 0.0;
    if (lane != 0) return;
    if (t.average_dump == p.dumpavg) {
        t.average_dump = 0;
        t.atlast = 0.0;
    }
    t.atlast = t.avgt;
    // distance test every dumpdist steps (cid.f90:997-1026)
    if (t.distance_dump == p.dumpdist) {
        t.distance_dump = 0;
        double cm[3];
        md_center_of_mass(nuc, mass, x, cm);
        const double d = cid_dist(t.nuc0, x, cm);
        if (d > t.highest) t.highest = d;
        if (d < t.lowest) t.lowest = d;
        if (t.lowest < d) t.step_counter = t.step_counter + 1;
        else t.step_counter = 0;
        if (t.step_counter == 5) {
            t.total_steps = t.nstep + (int)llround(800.0 * (2.0 * p.tstep * (1.0 / kFsToAu)));
            t.collided = 1;
            t.collision_step = t.nstep;
            t.tav = 0.0;
            t.m = 0;
        }
    }
    if (nfrag > 1 && t.collided && !t.fragmented) {
        t.total_steps = t.nstep + t.add_steps;
        t.fragmented = 1;
    }
    if (t.nstep >= t.total_steps) {
        t.velo_diff = t.velo_cm - t.new_velo;
        t.done = 1;
        C.active[sys] = 0;
        // fragment energies and temperatures (md.f90:715-741)
        int n[10];
        for (int f = 0; f < 10; ++f) {
            t.eint[f] = 0.0;
            n[f] = 0;
        }
        const double* v = C.velo + 3 * (size_t)off;
        for (int i = 0; i < nuc; ++i) {
            const int j = C.list[off + i];
            if (j < 1 || j > 10) continue;
            t.eint[j - 1] = t.eint[j - 1] + 0.5 * mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] +
                                                             v[3 * i + 2] * v[3 * i + 2]);
            n[j - 1] = n[j - 1] + 1;
        }
        for (int f = 0; f < nfrag && f < 10; ++f) t.fragT[f] = t.eint[f] / ((double)(1.5f * (float)n[f]) * kKB);
    }
}

}  // namespace qx
