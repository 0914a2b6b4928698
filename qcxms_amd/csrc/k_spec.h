// MO spectrum of one system after an SCC: orbital energies in ascending order,
// occupations and Mulliken MO populations per atom, as write_qmo
// (mo_energ.f90:7-80) derives them from tblite's wavefunction
// (tblite.f90:153-166) for getspec (mo_spec.f90:12-119).
#pragma once
#include "common.h"

namespace qx {

// One workgroup.  C = X U (AO coefficients; X = L^-T, U the eigenvectors in
// the orthogonalised basis) is formed in the system's T slot; orbital i goes
// to position rank(i) of the ascending order (ties by index).  For that
// position k and atom a:
//   qmo[k*nat + a] = (sum_{j on a} sum_m C_ji C_mi S_jm + 1e-10) / norm_k
// with norm_k the sum over atoms, as mo_energ.f90:36-59.  focc is the total
// occupation of the spatial orbital (2 f).
__global__ __launch_bounds__(QX_BLOCK) void k_mo_pop(BatchPtrs B, int sys, double* emo, double* focc,
                                                    double* qmo)
{
    __shared__ ShellDesc sh[QX_MAX_NSH];
    __shared__ int ao_at[QX_MAX_NAO], ao_sh[QX_MAX_NAO], rank[QX_MAX_NAO];
    const SysInfo si = B.info[sys];
    const int n = si.nao, ld = si.ld, nat = si.nat, tid = threadIdx.x;
    const double* X = B.X + si.mat_off;
    const double* U = B.U + si.mat_off;
    const double* S = B.S + si.mat_off;
    double* Cm = B.T + si.mat_off;
    const double* e = B.eps + (size_t)sys * QX_MAX_NAO;
    const double* f = B.focc + (size_t)sys * QX_MAX_NAO;

    build_shell_tables(si, B.elem, sh, ao_at, ao_sh);
    for (int i = tid; i < n; i += QX_BLOCK) {
        int r = 0;
        for (int j = 0; j < n; ++j) r += (e[j] < e[i]) || (e[j] == e[i] && j < i);
        rank[i] = r;
    }
    for (int idx = tid; idx < n * n; idx += QX_BLOCK) {
        const int mu = idx / n, i = idx - mu * n;
        double c = 0.0;
        for (int j = 0; j < n; ++j) c += X[(size_t)mu * ld + j] * U[(size_t)j * ld + i];
        Cm[(size_t)mu * ld + i] = c;
    }
    __syncthreads();
    for (int i = tid; i < n; i += QX_BLOCK) {
        const int k = rank[i];
        emo[k] = e[i];
        focc[k] = 2.0 * f[i];
        double* row = qmo + (size_t)k * nat;
        double acc = 0.0, norm = 0.0;
        for (int j = 0; j < n; ++j) {
            double sc = 0.0;
            for (int m = 0; m < n; ++m) sc += S[(size_t)j * ld + m] * Cm[(size_t)m * ld + i];
            acc += Cm[(size_t)j * ld + i] * sc;
            // AOs of an atom are contiguous: close the atom's sum at its last AO
            if (j == n - 1 || ao_at[j + 1] != ao_at[j]) {
                const double q = acc + 1e-10;
                row[ao_at[j]] = q;
                norm += q;
                acc = 0.0;
            }
        }
        for (int a = 0; a < nat; ++a) row[a] /= norm;
    }
}

}  // namespace qx
