// Post-MD fragment analysis of QCxMS for finished trajectories: fragment
// ionisation potentials from GFN2 single points (analyse, src/analyse.f90:
// 14-502, through eqm, src/iniqm.f90:241-451), Boltzmann charge assignment
// (boltz, src/utility.f90:469-498) and the qcxms.res records
// (manage_fragments, src/write_fragments.f90:12-457, EI branch).
//
// Host code.  The single points of every fragment of every trajectory (the
// neutral and each ion charge) go to the GPU as one ragged batch through
// qcxms_egrad_batch; what remains is per-trajectory bookkeeping.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace qx_post {

constexpr int kMaxFrag = 10;                      // fragm(10), natf(10) in write_fragments.f90
constexpr int kMaxCharge = 3;                     // |mchrg| handled here
constexpr int kMaxElem = 200;                     // fragel(200,10) in fragmass
constexpr double kAutoEV = 27.21138505;           // xtb_mctc_convert.f90:27
constexpr double kBoltzEh = 3.166808578545117e-06;  // kB, xtb_mctc_constants.f90:35
constexpr double kIpEtemp = 300.0;                // analyse.f90:270-273 (progi 8)
constexpr double kHuge = std::numeric_limits<double>::max();  // huge(0.0_wp)

// iniqm.f90:17-41
inline int core_electrons(int z)
{
    if (z <= 2) return 0;
    if (z <= 10) return 2;
    if (z <= 18) return 10;
    if (z <= 29) return 18;
    if (z <= 36) return 28;
    if (z <= 47) return 36;
    if (z <= 54) return 46;
    if (z <= 71) return 54;
    if (z <= 79) return 68;
    if (z <= 86) return 78;
    return 0;
}

// electrons_amount (iniqm.f90:45-63): valence electrons at total charge chrg
inline int valence_electrons(const int* z, int n, int chrg)
{
    int nel = 0;
    for (int i = 0; i < n; ++i) nel += z[i] - core_electrons(z[i]);
    return nel - chrg;
}

// eself (iniqm.f90:430-451), the energy eqm uses when a fragment has no
// electrons: sum_i z_i^2 gam_i / 2 + z_i^3 gam3_i / 3 with z_i the valence
// electron count, gam and gam3 from dftd4.f90:38-72 (REAL(4) literals
// widened to REAL(8)).  Tabulated for H, C, N and O; false for others.
inline bool eself(const int* z, int n, double& e)
{
    e = 0.0;
    for (int i = 0; i < n; ++i) {
        double g, g3;
        switch (z[i]) {
        case 1: g = (double)0.47259288f; g3 = (double)-0.02448f; break;
        case 6: g = (double)0.42195412f; g3 = (double)0.167160f; break;
        case 7: g = (double)0.50438193f; g3 = (double)0.156306f; break;
        case 8: g = (double)0.58691863f; g3 = (double)0.161466f; break;
        default: return false;
        }
        const double zv = (double)(z[i] - core_electrons(z[i]));
        e = e + 0.5 * (zv * zv) * g + (zv * zv * zv) * g3 / 3.0;
    }
    return true;
}

// boltz (utility.f90:469-498), units = 2 (eV): populations exp(-ip/f) / sum
// with f = temp kB autoev, over an nfrag x nchrg column-major block.
inline void boltz(int nfrag, int nchrg, double temp, const double* ip, double* out)
{
    const double f = temp * kBoltzEh * kAutoEV;
    double esum = 0.0;
    for (int i = 0; i < nfrag; ++i)
        for (int j = 0; j < nchrg; ++j) esum = esum + std::exp(-ip[i + j * nfrag] / f);
    for (int i = 0; i < nfrag; ++i)
        for (int j = 0; j < nchrg; ++j) out[i + j * nfrag] = std::exp(-ip[i + j * nfrag] / f) / esum;
}

// ---------------------------------------------------------------------------
// Fortran edit descriptors of the qcxms.res record: Fw.d and Iw, right
// justified, a field that does not fit is w asterisks.
inline void put_int(std::string& s, long long v, int w)
{
    char b[32];
    const int n = std::snprintf(b, sizeof b, "%*lld", w, v);
    s += (n > w) ? std::string(w, '*') : std::string(b);
}

inline void put_fixed(std::string& s, double v, int w, int d)
{
    char b[64];
    const int n = std::snprintf(b, sizeof b, "%*.*f", w, d, v);
    s += (n > w || !std::isfinite(v)) ? std::string(w, '*') : std::string(b);
}

// One record of write_fragments.f90:402-441 with the format
// (F10.7,i3,2i5,2i2,2x,i3,2x,20(i4,i3)) over the items
//   EI  (no collision number): q, mchrg, itrj, isec, j, l, (Z_m, n_m, m = 1..l)
//   CID (icoll >= 0)         : q, mchrg, itrj, icoll, isec, j, l, (Z_m, n_m, ...)
// Items and edit descriptors are paired in order, so the EI record puts j and
// l in the two i2 fields and the first Z in the i3 field; output ends at the
// first data descriptor left without an item.
inline std::string res_record(double q, int mchrg, int itrj, int icoll, int isec, int j, int l,
                              const int* zel, const int* cnt)
{
    std::vector<long long> items = {mchrg, itrj};
    if (icoll >= 0) items.push_back(icoll);
    items.push_back(isec);
    items.push_back(j);
    items.push_back(l);
    for (int m = 0; m < l; ++m) {
        items.push_back(zel[m]);
        items.push_back(cnt[m]);
    }
    std::string s;
    put_fixed(s, q, 10, 7);
    // descriptor list after F10.7; 0 stands for the 2x position edits
    static const int desc[] = {3, 5, 5, 2, 2, 0, 3, 0};
    size_t k = 0;
    for (int w : desc) {
        if (w == 0) {
            s += "  ";
            continue;
        }
        if (k == items.size()) return s;
        put_int(s, items[k++], w);
    }
    for (int rep = 0; rep < 20 && k < items.size(); ++rep) {
        put_int(s, items[k++], 4);
        if (k == items.size()) break;
        put_int(s, items[k++], 3);
    }
    return s;
}

// ---------------------------------------------------------------------------
// One trajectory's fragment bookkeeping after its IPs are known: the
// charge-assignment part of manage_fragments (write_fragments.f90:112-300).
struct ChargeState {
    int nfrag = 0;
    int mchrg = 1;          // in/out
    double chrgcont = 1.0;  // in/out
    int tcont = 0;          // out: fragment the run continues with (0: none)
    double fragchrg3[kMaxFrag] = {0};
};

// ip: fragip(nfrag, |mchrg|) column-major (eV); natf / first element per
// fragment for the single-H-atom rule.  temp = aTlast * btf.
inline void assign_charges(ChargeState& cs, const double* fragip, const int* natf, const int* zfirst,
                           double temp, int isec, int maxsec)
{
    const int nfrag = cs.nfrag, nch = std::abs(cs.mchrg);
    double q2[kMaxFrag * kMaxCharge] = {0};
    if (nfrag > 1) {
        // relative ionisation potentials 0 -> 1, 1 -> 2, ... (write_fragments.f90:114-135)
        double ipd[kMaxFrag * kMaxCharge];
        for (int i = 0; i < nfrag; ++i)
            for (int j = 1; j <= nch; ++j) {
                const double prev = (j == 1) ? 0.0 : fragip[i + (j - 2) * nfrag];
                double d = fragip[i + (j - 1) * nfrag] - prev;
                if (natf[i] == 1 && zfirst[i] == 1) d = (j == 1) ? fragip[i] : kHuge;
                ipd[i + (j - 1) * nfrag] = d;
            }
        if (nch > 1) {
            // rank the IP differences and Boltzmann-weight only the charges
            // beyond the first nch - 1 (write_fragments.f90:152-218)
            const int n = nfrag * nch;
            int pos[kMaxFrag * kMaxCharge];
            int nr = 0;
            double last = -kHuge;
            for (int c = 0; c < n; ++c) {
                int best = -1;
                for (int k = 0; k < n; ++k)
                    if (ipd[k] > last && (best < 0 || ipd[k] < ipd[best])) best = k;
                if (best < 0) break;
                pos[nr++] = best;
                last = ipd[best];
            }
            double ipd2[kMaxFrag * kMaxCharge];
            for (int k = 0; k < n; ++k) ipd2[k] = kHuge;
            if (nr >= nch)
                for (int c = nch - 1; c < nr; ++c) ipd2[pos[c]] = ipd[pos[c]] - ipd[pos[nch - 2]];
            for (int c = 0; c < nch - 1 && c < nr; ++c) ipd2[pos[c]] = kHuge;
            boltz(nfrag, nch, temp, ipd2, q2);
            for (int c = 0; c < nch - 1 && c < nr; ++c) q2[pos[c]] = 1.0;
        } else {
            boltz(nfrag, nch, temp, ipd, q2);
        }
    }
    if (cs.mchrg < 0)
        for (int k = 0; k < nfrag * nch; ++k) q2[k] = -q2[k];
    for (int i = 0; i < kMaxFrag; ++i) cs.fragchrg3[i] = 0.0;
    if (nfrag > 1) {
        for (int i = 0; i < nfrag; ++i)
            for (int j = 0; j < nch; ++j) cs.fragchrg3[i] = cs.fragchrg3[i] + q2[i + j * nfrag];
    } else {
        for (int i = 0; i < kMaxFrag; ++i) cs.fragchrg3[i] = cs.chrgcont;
    }
    // fragment the trajectory continues with (write_fragments.f90:249-272)
    cs.tcont = 0;
    if (cs.mchrg > 0) {
        double largest = -1.0;
        for (int i = 0; i < nfrag; ++i) {
            const double w = (double)(float)natf[i] * cs.fragchrg3[i];
            if (w > largest) {
                largest = w;
                cs.tcont = i + 1;
            }
        }
    } else {
        double largest = 1.0;
        for (int i = 0; i < nfrag; ++i) {
            const double w = (double)(float)natf[i] * cs.fragchrg3[i];
            if (w < largest) {
                largest = w;
                cs.tcont = i + 1;
            }
        }
    }
    // fractional charges and the charge carried on (write_fragments.f90:278-311)
    if (nfrag == 1 || isec == maxsec + 1) {
        cs.tcont = 0;
    } else {
        for (int i = 0; i < nfrag; ++i)
            cs.fragchrg3[i] = cs.fragchrg3[i] * std::fabs(cs.chrgcont) / std::abs(cs.mchrg);
        if (cs.tcont > 0) {
            cs.chrgcont = cs.fragchrg3[cs.tcont - 1];
            const int nearest = (int)std::lround(cs.fragchrg3[cs.tcont - 1]);
            if (nearest > 0 && cs.mchrg > 0) cs.mchrg = nearest;
            if (nearest == 0 && cs.mchrg > 0) cs.mchrg = 1;
            if (nearest < 0 && cs.mchrg < 0) cs.mchrg = nearest;
            if (nearest == 0 && cs.mchrg < 0) cs.mchrg = -1;
        }
    }
}

}  // namespace qx_post
